/*
 * bms2.h -- C ABI of the MI355X (gfx950) batched subgraph-isomorphism library.
 *
 * Every entry point returns an int status (BMS2_OK == 0); bms2_last_error() gives the message of
 * the last failure on the calling thread.  Device work is enqueued on a bms2_queue (a HIP stream);
 * calls are ordered on that queue and return before the work completes unless stated otherwise.
 * Plain pointers and sizes only: no C++ or torch types cross this boundary.
 *
 * Each function names the reference interface it replaces (paths relative to the AnySparse/BMS2
 * tree).  The C++ facade in include/sigmo.hpp wraps these calls behind the reference's own class
 * and function names.
 */
#ifndef BMS2_H
#define BMS2_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status codes */
#define BMS2_OK 0
#define BMS2_ERR_INVALID 1     /* bad argument, shape or handle */
#define BMS2_ERR_UNSUPPORTED 2 /* input outside the supported limits below */
#define BMS2_ERR_HIP 3         /* HIP runtime failure */
#define BMS2_ERR_IO 4          /* file could not be opened / written */
#define BMS2_ERR_FORMAT 5      /* malformed .dat line (graph.hpp:199-201) */
#define BMS2_ERR_NOMEM 6       /* host or device allocation failed */
#define BMS2_ERR_CAPACITY 7    /* caller buffer too small */

/* ---------------------------------------------------------------- constants
 * types.hpp:18 WILDCARD_NODE; isomorphism.hpp:248 MAX_QUERY_NODES (=30 there, 32 here);
 * data graphs up to 256 nodes (the reference is exact only up to 64, utils.hpp:95-97). */
#define BMS2_WILDCARD_NODE 255u
#define BMS2_MAX_QUERY_NODES 32u
#define BMS2_MAX_GRAPH_NODES 256u
#define BMS2_SCOPE_DATA 0
#define BMS2_SCOPE_QUERY 1

const char* bms2_last_error(void);
/* Library version as 10000*major + 100*minor + patch. */
int bms2_version(void);

/* ---------------------------------------------------------------- host containers
 * A packed batch of CSR graphs in host memory, the layout createDeviceCSRGraph packs
 * (graph.hpp:251-290): graph_offsets is the node prefix over graphs, column_indices hold global
 * node ids, each undirected edge appears under both endpoints in file edge order
 * (graph.hpp:229-240).  Arrays filled by the library are released with bms2_host_csr_free. */
typedef struct bms2_host_csr {
  uint32_t num_graphs;
  uint32_t reserved;
  uint64_t total_nodes;
  uint64_t total_edges; /* directed CSR entries */
  uint32_t* graph_offsets;  /* [num_graphs + 1] */
  uint32_t* row_offsets;    /* [total_nodes + 1] */
  uint32_t* column_indices; /* [total_edges] */
  uint8_t* node_labels;     /* [total_nodes] */
  uint8_t* edge_labels;     /* [total_edges] */
} bms2_host_csr;

void bms2_host_csr_free(bms2_host_csr* g);

/* Reads a .dat file, one graph per line: "n#N l#L (id label)xN e#E (u v [label])xE".
 * Replaces io::loadCSRGraphsFromFile (io.hpp:43-52) with IntermediateGraph (graph.hpp:150-202)
 * and toCSRGraph (graph.hpp:212-243).  max_graphs == 0 reads every line.  A malformed line
 * fails with BMS2_ERR_FORMAT naming the line, unless skip_malformed is set, in which case it is
 * dropped as BMS2_batch.cpp:121-124 does. */
int bms2_load_dat(const char* path, uint64_t max_graphs, int skip_malformed, bms2_host_csr* out);
/* Writes the batch in the same format (the writer of generate_data.py:76-92). */
int bms2_write_dat(const char* path, const bms2_host_csr* g);
/* io::loadCSRGraphsFromLines (io.hpp:34-41): the same reader over num_lines lines in memory, one
 * graph per line (the batch driver's ingestion, BMS2_batch.cpp:104-124).  skip_malformed drops
 * empty and malformed lines, as the batch driver's try/catch does; otherwise the first one fails
 * with BMS2_ERR_FORMAT naming its index. */
int bms2_parse_dat_lines(const char* const* lines, uint64_t num_lines, int skip_malformed,
                         bms2_host_csr* out);

/* Deterministic synthetic molecule workloads (no reference counterpart: the reference reads
 * RDKit-converted SMARTS files, generate_data.py).  Graph i of the data batch depends only on
 * (seed, i), so any prefix of a large batch equals the smaller batch. */
typedef struct bms2_workload_spec {
  uint64_t seed;
  uint32_t num_data_graphs;
  uint32_t num_query_graphs;
  uint32_t data_size_dist;       /* 0: uniform [min,max]; 1: lognormal(mu, sigma) clipped */
  uint32_t data_min_nodes;
  uint32_t data_max_nodes;
  uint32_t polycyclic_permille;  /* share of data graphs built as fused ring systems */
  double data_lognormal_mu;
  double data_lognormal_sigma;
  uint32_t poly_min_rings;
  uint32_t poly_max_rings;
  uint32_t query_min_nodes;
  uint32_t query_max_nodes;
  uint32_t extracted_permille;   /* share of queries cut from data graphs (rest independent) */
  uint32_t query_shapes;         /* 0: molecule-like; 1: adds ring and branched shapes */
  uint32_t max_ring_closures;    /* ring-closing edges per molecule: uniform 0..max */
  uint32_t num_threads;          /* 0: hardware concurrency */
} bms2_workload_spec;

int bms2_generate_workload(const bms2_workload_spec* spec, bms2_host_csr* data_out,
                           bms2_host_csr* query_out);

/* ---------------------------------------------------------------- queues and events */
typedef struct bms2_queue_s* bms2_queue;
typedef struct bms2_event_s* bms2_event;

/* A queue owns a new HIP stream on `device` (sycl::queue in BMS2.cpp:14). */
int bms2_queue_create(int device, bms2_queue* out);
/* A queue over an existing hipStream_t (not owned). */
int bms2_queue_wrap(int device, void* hip_stream, bms2_queue* out);
/* Releases the queue (and its stream when owned).  A queue keeps a grow-only scratch buffer for
 * the filter, refinement and join calls until it is destroyed, sized for the largest batch they
 * ran on plus an eighth, at least 1 MiB: the filter's 64 B per query node, the refinement's 8 B
 * per data node, the join's 8 B per data node + 128 B per query graph (graphs of up to 64 nodes)
 * or, above 64 nodes, 16 B per data node (32 B above 128 nodes) + 128 B per query graph + 8 B
 * per GMCR pair + the join lanes' stashes (16 B per query depth and row word for each of 12
 * waves per CU: about 100 MB for queries of up to 16 nodes and rows of 2 words).
 * A queue is used from one host thread at a time (like one stream's launches, calls from two
 * threads would interleave and share the scratch buffer); use a queue per thread. */
int bms2_queue_destroy(bms2_queue q);
int bms2_queue_wait(bms2_queue q); /* queue.wait_and_throw() */
void* bms2_queue_stream(bms2_queue q);

/* Every stage call takes an optional bms2_event* out: when non-null it receives a new event
 * pair bracketing the call's kernels (utils::BatchedEvent, utils.hpp:123-157). */
int bms2_event_elapsed_ms(bms2_event e, float* ms); /* waits for the event */
int bms2_event_destroy(bms2_event e);

/* The queue's device: global memory bytes, compute units and name (name_capacity bytes, NUL
 * terminated); every output is nullable.  Replaces queue.get_device().get_info<...>() of
 * BMS2.cpp:15-16 and device::getDeviceMemorySize (device.hpp:19-21). */
int bms2_queue_device_info(bms2_queue q, uint64_t* global_mem_bytes, uint32_t* compute_units,
                           char* name, uint64_t name_capacity);

/* Device memory helpers for callers without another allocator. */
int bms2_device_malloc(bms2_queue q, uint64_t bytes, void** out);
/* Pinned host memory that kernels read and write directly (sycl::malloc_shared in BMS2.cpp:205,
 * 243-244, 269): the drivers' match totals, partial-match buffers and pair flags.  Released with
 * bms2_shared_free, which first waits for the queue. */
int bms2_shared_malloc(bms2_queue q, uint64_t bytes, void** out);
int bms2_shared_free(bms2_queue q, void* ptr);
int bms2_device_free(bms2_queue q, void* ptr);
int bms2_memcpy_h2d(bms2_queue q, void* dst, const void* src, uint64_t bytes);
int bms2_memcpy_d2h(bms2_queue q, void* dst, const void* src, uint64_t bytes); /* synchronous */
int bms2_memset_zero(bms2_queue q, void* dst, uint64_t bytes);

/* ---------------------------------------------------------------- device graphs
 * Same fields as sigmo::DeviceBatchedCSRGraph (graph.hpp:58-66), device pointers. */
typedef struct bms2_device_csr {
  uint32_t* graph_offsets;
  uint32_t* row_offsets;
  uint32_t* column_indices;
  uint8_t* node_labels;
  uint8_t* edge_labels;
  uint32_t num_graphs;
  uint32_t max_graph_nodes; /* largest graph, filled at creation */
  uint64_t total_nodes;
  uint64_t total_edges;
  uint32_t max_degree; /* largest CSR row, filled at creation */
  uint32_t reserved;
} bms2_device_csr;

/* createDeviceCSRGraph (graph.hpp:251-310): uploads a packed host batch. */
int bms2_device_csr_create(bms2_queue q, const bms2_host_csr* host, bms2_device_csr* out);
/* createDeviceCSRGraph with the packing done on the GPU (the reference packs on the host in a
 * serial loop marked "TODO offload on GPU", graph.hpp:249-310).  The graphs arrive as their own
 * local arrays laid end to end, with no offsets added: nodes[g] node counts; ro_local[total_nodes]
 * for each node the end of its neighbour list counted from its graph's first CSR entry (a
 * CSRGraph's row_offsets[1..n]); ci_local[total_edges] neighbour ids local to the graph;
 * node_labels; edge_labels (nullable: all 0).  The device turns them into the global batch and
 * checks what bms2_device_csr_create checks on the host (endpoints inside their graph, no
 * self-loops, symmetric lists, graphs within BMS2_MAX_GRAPH_NODES). */
int bms2_device_csr_pack(bms2_queue q, uint32_t num_graphs, const uint32_t* nodes,
                         const uint32_t* ro_local, const uint32_t* ci_local,
                         const uint8_t* node_labels, const uint8_t* edge_labels,
                         uint64_t total_nodes, uint64_t total_edges, bms2_device_csr* out);
/* destroyDeviceCSRGraph (graph.hpp:312-318). */
int bms2_device_csr_destroy(bms2_queue q, bms2_device_csr* g);
/* getDeviceCSRGraphAllocSize (graph.hpp:320-324). */
uint64_t bms2_device_csr_alloc_size(const bms2_device_csr* g);

/* ---------------------------------------------------------------- Signature<PowerGraph, 4>
 * 16 x 4-bit saturating label histograms (signature.hpp:22-53).  Host copies are one u64 per
 * node, the reference's SignatureDevice layout. */
typedef struct bms2_signature_s* bms2_signature;
/* Signature(queue, data_nodes, query_nodes) (signature.hpp:156-167); reachability state is
 * zero-initialised. */
int bms2_signature_create(bms2_queue q, uint64_t data_nodes, uint64_t query_nodes,
                          bms2_signature* out);
int bms2_signature_destroy(bms2_signature s);
/* getDataSignatureAllocationSize / getQuerySignatureAllocationSize (signature.hpp:180-193): device
 * bytes held for one scope (40 per node: the histogram and four u64 planes of counted nodes). */
uint64_t bms2_signature_alloc_size(bms2_signature s, int scope);
/* generateDataSignatures / generateQuerySignatures (signature.hpp:56-84, 124-151). */
int bms2_signature_generate(bms2_signature s, const bms2_device_csr* g, int scope, bms2_event* ev);
/* refineDataSignatures / refineQuerySignatures (signature.hpp:61-73, 336-389): histogram of the
 * nodes within view_size hops, minus the node and every node an earlier call already counted.
 * As in the reference, those counted sets persist for the object's lifetime (generate does not
 * clear them): each run of the stage sequence needs a new Signature, or its refinements come out
 * empty, refine nothing and leave a larger GMCR than the previous run's. */
int bms2_signature_refine(bms2_signature s, const bms2_device_csr* g, uint32_t view_size, int scope,
                          bms2_event* ev);
int bms2_signature_copy_to_host(bms2_signature s, int scope, uint64_t* out);
const uint64_t* bms2_signature_device_ptr(bms2_signature s, int scope);

/* ---------------------------------------------------------------- PathSignature
 * 256-bit Bloom filter of label paths of 2 and 3 edges (signature.hpp:406-532).  Host copies are
 * four u64 per node ([node][4], the reference's PathSignatureDevice layout). */
typedef struct bms2_path_signature_s* bms2_path_signature;
int bms2_path_signature_create(bms2_queue q, uint64_t data_nodes, uint64_t query_nodes,
                               bms2_path_signature* out); /* signature.hpp:440-447 */
int bms2_path_signature_destroy(bms2_path_signature s);
/* signature.hpp:521-522: 32 bytes per node of the scope. */
uint64_t bms2_path_signature_alloc_size(bms2_path_signature s, int scope);
int bms2_path_signature_generate(bms2_path_signature s, const bms2_device_csr* g, int scope,
                                 bms2_event* ev); /* signature.hpp:455-519 */
int bms2_path_signature_copy_to_host(bms2_path_signature s, int scope, uint64_t* out);

/* ---------------------------------------------------------------- CycleSignature
 * Per node counts of 5- and 6-rings, same-label vs mixed (signature.hpp:535-732).  Host copies
 * are one u64 per node: u16 fields {5 same, 5 diff, 6 same, 6 diff} from the low end. */
typedef struct bms2_cycle_signature_s* bms2_cycle_signature;
int bms2_cycle_signature_create(bms2_queue q, uint64_t data_nodes, uint64_t query_nodes,
                                bms2_cycle_signature* out); /* signature.hpp:557-566 */
int bms2_cycle_signature_destroy(bms2_cycle_signature s);
/* signature.hpp:721-722: 8 bytes per node of the scope. */
uint64_t bms2_cycle_signature_alloc_size(bms2_cycle_signature s, int scope);
int bms2_cycle_signature_generate(bms2_cycle_signature s, const bms2_device_csr* g, int scope,
                                  bms2_event* ev); /* signature.hpp:574-719 */
int bms2_cycle_signature_copy_to_host(bms2_cycle_signature s, int scope, uint64_t* out);

/* ---------------------------------------------------------------- Candidates
 * Dense domain bit-matrix: one row per global query node, bit d%64 of u64 word d/64 for global
 * data node d (candidates.hpp:12-159, whose u32 words hold the same bits). */
typedef struct bms2_candidates_s* bms2_candidates;
int bms2_candidates_create(bms2_queue q, uint64_t source_nodes, uint64_t target_nodes,
                           bms2_candidates* out); /* candidates.hpp:161-175, zero-filled */
int bms2_candidates_destroy(bms2_candidates c);
uint64_t bms2_candidates_words_per_row(bms2_candidates c);
int bms2_candidates_copy_to_host(bms2_candidates c, uint64_t* out); /* candidates.hpp:181-186 */
/* Per-row popcounts (getCandidatesCount(source_node), candidates.hpp:73-80) into out[rows]. */
int bms2_candidates_row_counts(bms2_candidates c, uint32_t* out);
/* Rows [first_row, first_row + num_rows) of the matrix (num_rows x words_per_row u64) into out:
 * what getCandidatesCount(source_node[, start, end]) and getCandidateAt read on the host
 * (candidates.hpp:73-158). */
int bms2_candidates_copy_rows(bms2_candidates c, uint64_t first_row, uint64_t num_rows,
                              uint64_t* out);

/* ---------------------------------------------------------------- filter / refine
 * filterCandidates (isomorphism.hpp:30-104): sets bit (q, d) when the labels agree (or the query
 * label is the wildcard), deg(q) <= deg(d), path(d) contains path(q) and every cycle count of d
 * is >= that of q.  Bits are OR-ed into the matrix. */
int bms2_filter_candidates(bms2_queue q, const bms2_device_csr* query, const bms2_device_csr* data,
                           bms2_signature sig, bms2_path_signature path,
                           bms2_cycle_signature cycle, bms2_candidates cand, bms2_event* ev);
/* refineCandidates (isomorphism.hpp:106-203): clears (q, d) unless every label count of the
 * current query signature is <= that of the data signature. */
int bms2_refine_candidates(bms2_queue q, const bms2_device_csr* query, const bms2_device_csr* data,
                           bms2_signature sig, bms2_candidates cand, bms2_event* ev);

/* ---------------------------------------------------------------- GMCR
 * mapping::GMCR (gmcr.hpp:16-145): CSR from data graph to the query graphs with more than one
 * node whose every node keeps a candidate in that data graph; ascending query id per segment.
 * Offsets are u32 as in the reference: a batch with more than 2^32 - 1 GMCR pairs fails with
 * BMS2_ERR_UNSUPPORTED. */
typedef struct bms2_gmcr_s* bms2_gmcr;
int bms2_gmcr_create(bms2_queue q, bms2_gmcr* out); /* gmcr.hpp:26 */
int bms2_gmcr_destroy(bms2_gmcr g);
int bms2_gmcr_generate(bms2_gmcr g, const bms2_device_csr* query, const bms2_device_csr* data,
                       bms2_candidates cand, bms2_event* ev); /* gmcr.hpp:35-142 */
uint64_t bms2_gmcr_total(bms2_gmcr g);
int bms2_gmcr_device_ptrs(bms2_gmcr g, const uint32_t** data_graph_offsets,
                          const uint32_t** query_graph_indices);
int bms2_gmcr_copy_to_host(bms2_gmcr g, uint32_t* data_graph_offsets /*[D+1]*/,
                           uint32_t* query_graph_indices /*[total]*/);

/* ---------------------------------------------------------------- join
 * joinCandidates (isomorphism.hpp:291-472): for every GMCR pair, the number of injective maps f
 * with f(i) in the domain of query node i inside the data graph and query adjacency equal to data
 * adjacency on every mapped pair (induced).  find_first counts 1 per pair with any embedding.
 * d_num_matches is a device u64 the call atomically adds the total to (the caller zeroes it, as
 * BMS2.cpp:205-206 does).  d_pair_counts (nullable) receives one u64 per GMCR entry; it holds
 * pair_capacity entries, and a call whose GMCR has more (bms2_gmcr_total) fails with
 * BMS2_ERR_CAPACITY before any device work is enqueued. */
int bms2_join_candidates(bms2_queue q, const bms2_device_csr* query, const bms2_device_csr* data,
                         bms2_candidates cand, bms2_gmcr gmcr, uint64_t* d_num_matches,
                         int find_first, uint64_t* d_pair_counts, uint64_t pair_capacity,
                         bms2_event* ev);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): bms2_join_candidates
 * with no event and no allocation of any kind.  The join's scratch (adjacency rows of the data
 * nodes, the queries' depth adjacency, work cursors: 8 bytes per data node + 128 per query graph
 * when every data graph has at most 64 nodes; otherwise 16 bytes per data node (32 above 128
 * nodes) + 128 per query graph + 8 per GMCR pair) lives in a static device array of
 * BMS2_JOIN_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until the process
 * ends; a batch that needs more fails with BMS2_ERR_CAPACITY.  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (from two streams or two host threads) race. */
#define BMS2_JOIN_NOALLOC_SCRATCH_BYTES (64ull << 20)
int bms2_join_candidates_noalloc(bms2_queue q, const bms2_device_csr* query,
                                 const bms2_device_csr* data, bms2_candidates cand, bms2_gmcr gmcr,
                                 uint64_t* d_num_matches, int find_first, uint64_t* d_pair_counts,
                                 uint64_t pair_capacity);

/* CS (cluster) join, step 1: joinPartialCandidates2 (isomorphism.hpp:1021-1178), the --use-cs
 * path of BMS2.cpp:217-259.  Every embedding of query 0's first `end` nodes (the cluster
 * prototype; BMS2 --use-cs takes end = its node count) in every data graph, each node inside its
 * domain, injective, induced adjacency at every depth, standard backtracking.  Written to d_partial
 * as records of end + 1 int32 {mapped global data node ids in depth order, data graph id}, in data
 * graph order then ascending DFS order (the reference's order depends on atomics).  *num_partial
 * (host) receives the number found; when it exceeds `capacity` (records) nothing is written and
 * the call fails with BMS2_ERR_CAPACITY (the reference writes past its 10^8-record buffer).
 * Synchronises the queue. */
int bms2_join_partial_candidates(bms2_queue q, const bms2_device_csr* query,
                                 const bms2_device_csr* data, bms2_candidates cand, uint32_t end,
                                 int32_t* d_partial, uint64_t capacity, uint64_t* num_partial,
                                 bms2_event* ev);

/* CS (cluster) join, step 2: joinWithPartialMatches (isomorphism.hpp:756-977).  For every partial
 * match of step 1 and every GMCR query q of its data graph: n_q == end counts the match itself,
 * n_q > end counts its extensions to depths end..n_q-1 (q's domains, q's induced adjacency against
 * every earlier depth including the prototype part, injective), n_q < end nothing.
 * d_pair_counts (required, one u64 per GMCR entry, zeroed by the call) receives the counts, or in
 * find_first mode a 0/1 flag per entry (the reference's pair_done flags); it holds pair_capacity
 * entries, and a GMCR with more fails with BMS2_ERR_CAPACITY before any device work.  The total
 * is atomically added to the device u64 *d_num_matches (the caller zeroes it). */
int bms2_join_with_partial_matches(bms2_queue q, const bms2_device_csr* query,
                                   const bms2_device_csr* data, bms2_candidates cand,
                                   bms2_gmcr gmcr, uint32_t end, const int32_t* d_partial,
                                   uint64_t num_partial, uint64_t* d_num_matches, int find_first,
                                   uint64_t* d_pair_counts, uint64_t pair_capacity,
                                   bms2_event* ev);

/* ---------------------------------------------------------------- fused pipeline
 * The stage sequence of BMS2.cpp:131-234 (signatures -> filter -> refine x K -> GMCR -> join) as
 * fused launches that keep domains on chip.  Results equal the staged calls above. */
typedef struct bms2_pipeline_options {
  uint32_t refinement_steps; /* K (BMS2 -i) */
  uint32_t find_first;       /* 1: first-match, 0: count-all (BMS2 --find-all) */
  uint32_t block_waves;      /* waves per CU of the small-query join; 0 = tuned */
  uint32_t join_waves;       /* waves per CU of the large-query join; 0 = tuned */
  uint32_t no_graph;         /* 1: never capture the chunk kernels in a hipGraph */
  uint32_t reserved[3];
} bms2_pipeline_options;

typedef struct bms2_pipeline_s* bms2_pipeline;
int bms2_pipeline_create(bms2_queue q, const bms2_device_csr* query, const bms2_device_csr* data,
                         const bms2_pipeline_options* opt, bms2_pipeline* out);
int bms2_pipeline_destroy(bms2_pipeline p);
/* Runs the whole sequence.  Once a run has measured how many pair records the batch produces
 * and they fit one chunk, the chunk's kernels are captured in a hipGraph (per output buffers
 * and K) and later runs replay it; a replay that overflows drops the graph and redoes the run.
 * d_counts: u64[num_data_graphs * num_query_graphs], data-graph major
 * (pair (g, q) at g*Q + q), 0 for pairs outside the GMCR.  d_valid: u64 bit rows
 * [num_data_graphs][ceil(Q/64)], bit q set when (q, g) is a GMCR pair.  d_stats (nullable)
 * u64[8]: {GMCR pairs, matched pairs, total embeddings, DFS nodes expanded, 0...}.  ev brackets
 * the whole call; kernel_ev (nullable) only the fused screen + join kernel. */
int bms2_pipeline_run(bms2_pipeline p, uint64_t* d_counts, uint64_t* d_valid, uint64_t* d_stats,
                      bms2_event* ev, bms2_event* kernel_ev);

/* Binds another query batch to the pipeline's resident data batch (the data plan, its packs and
 * adjacency rows are kept; the matching order returns to identity, the launch shape of the
 * small-query join is kept).  The reference regenerates query signatures per data batch
 * (BMS2_batch.cpp:162-283); here the data side is the resident part.  The caller keeps the query
 * CSR alive until the next bind or destroy. */
int bms2_pipeline_set_queries(bms2_pipeline p, const bms2_device_csr* query);

/* Per query of the bound batch, over d_counts of the last run (u64[D*Q]): the number of data
 * graphs with a non-zero count (d_matched u64[Q]) and the summed counts (d_embeddings u64[Q]).
 * Serving callers read these instead of the dense counts.  No counterpart in the reference. */
int bms2_pipeline_query_totals(bms2_pipeline p, const uint64_t* d_counts, uint64_t* d_matched,
                               uint64_t* d_embeddings);

/* Launch shape of the joins, picked by bms2_pipeline_autotune. */
typedef struct bms2_pipeline_tuning {
  uint32_t small_waves; /* waves per CU of the small-query join */
  uint32_t large_waves; /* waves per CU of the large-query join */
  uint32_t slots;       /* pack-adjacency slots per small-join wave (8 or 16) */
  uint32_t candidates;  /* configurations timed */
  float best_ms;        /* fastest kernel span of one run */
  /* joins of the data graphs above 64 nodes (tuned when the batch has any) */
  uint32_t big_waves;       /* waves per CU */
  uint32_t big_block_waves; /* waves per workgroup: 1, 2 or 4 */
  uint32_t big_row_words;   /* u64 words per adjacency row of graphs of 65-128 nodes: 2 or 4 */
} bms2_pipeline_tuning;

/* Times whole runs of the resident batch under a set of launch shapes (waves per CU of each join,
 * pack-adjacency slots per small-join wave; for batches with graphs above 64 nodes also the waves
 * per CU, waves per workgroup and adjacency-row words of their joins), `runs` timed runs each
 * after one untimed run, and keeps the fastest for later runs.  The outputs hold the last run's results (identical under
 * every shape).  Refinement depth is not tuned: it changes the reference's results (shells at
 * distance >= 3 can remove true embeddings).  No counterpart in the reference. */
int bms2_pipeline_autotune(bms2_pipeline p, uint64_t* d_counts, uint64_t* d_valid, uint32_t runs,
                           bms2_pipeline_tuning* out);

/* Matching order of the join, per query: order[q*32 + d] = local query node matched at DFS
 * depth d (a permutation of 0..n_q-1 per query; NULL restores identity, the reference's order,
 * isomorphism.hpp:371-447).  Counts do not depend on the order (standard backtracking); only the
 * number of DFS nodes does, and an order that matches a node before any of its neighbours leaves
 * that depth unconstrained, which on large queries can grow the search combinatorially.  The
 * reference's join fixes identity order; defineMatchingOrder (isomorphism.hpp:476-502) is its
 * unused alternative. */
int bms2_pipeline_set_order(bms2_pipeline p, const uint8_t* order);

/* Isomorphic queries of one batch have equal results against every data graph: each node's
 * filter and refinement fields (label, degree, path Bloom filter, ring counts, shell histograms,
 * isomorphism.hpp:71-97,142-196) depend on labels and adjacency only, so an isomorphism carries
 * domains to domains and embeddings to embeddings.  bms2_pipeline_run therefore screens and joins
 * the lowest-indexed query of each isomorphism class (bms2_query_classes, computed when the batch
 * is bound) and, after the joins, copies its count and GMCR bit to the class's other queries in
 * every data graph (inside the timed kernel span).  d_counts, d_valid and stats[0..2] are those of
 * screening every query; stats[3] counts the DFS nodes actually expanded (also reduced by the
 * large-query join's symmetry breaking, bms2_query_symmetry).  on = 0 screens every query and
 * searches every embedding (traces always do).  *num_dups (nullable) = queries that copy their class's results.  No
 * counterpart in the reference, which runs every query of a batch. */
int bms2_pipeline_set_dedup(bms2_pipeline p, int on, uint32_t* num_dups);

/* Host: rep[g] = lowest-indexed query of the batch isomorphic to query g (labels and induced
 * adjacency preserved; edge labels ignored, as the join ignores them), each class proven by an
 * explicit isomorphism; queries with parallel edges, self-loops or above 32 nodes, and searches
 * over budget, stay alone.  *num_dups (nullable) = queries with rep[g] != g. */
int bms2_query_classes(const bms2_host_csr* q, uint32_t* rep, uint32_t* num_dups);

/* Host: orbit[v] (v a global node id of the batch) = the lowest local node of v's orbit under its
 * query's label-preserving automorphisms, each orbit proven by explicit automorphisms (a node
 * stays alone when a search runs over budget or its query has parallel edges or self-loops).
 * Automorphic nodes have equal filter and refinement fields, hence equal domains: the fused
 * screen tests one node per orbit of a query of up to 8 nodes and copies its domain to the rest. */
int bms2_query_orbits(const bms2_host_csr* q, uint8_t* orbit);

/* Host: the symmetry breaking the fused pipeline's count-all joins apply on data graphs of up to
 * 64 nodes (with class copies on, bms2_pipeline_set_dedup) under a matching order (order[q*32 + d] = local node
 * at depth d; NULL = identity).  Per query, of its automorphism orbits of two or more nodes the
 * one whose second-matched member is earliest: du[q] = the depth of its first-matched member,
 * vmask[q] = the depths of the others (du[q] = 0xff, vmask[q] = 0: none).  The join takes at a
 * depth of vmask only candidates above the data node matched at depth du and multiplies the
 * count by popcount(vmask) + 1: automorphic nodes have equal domains, so the automorphisms permute
 * a pair's embeddings freely and exactly one in every |orbit| puts the orbit's smallest data node
 * at depth du.  No counterpart in the reference. */
int bms2_query_symmetry(const bms2_host_csr* q, const uint8_t* order, uint8_t* du, uint32_t* vmask);

/* Data packs are planned in windows of BMS2_PACK_WINDOW consecutive data graphs. */
#define BMS2_PACK_WINDOW 1024u

/* Join traces over the data graphs of the leading pack windows: the pipeline runs on graphs
 * [0, *graphs_out), where *graphs_out is max_graphs rounded up to a multiple of BMS2_PACK_WINDOW
 * (all graphs when the batch is smaller), and accumulates, per query
 * node (global query node id), the summed domain size over its GMCR pairs (d_domsum u64[Qn]),
 * per query the number of GMCR pairs (d_qpairs u64[Q]) and DFS nodes expanded (d_qnodes u64[Q]),
 * and per pair the DFS nodes (d_pair_nodes u32[graphs][Q], saturating, nullable).  The caller
 * zeroes the accumulators.  d_counts / d_valid are sized for the traced graphs as in
 * bms2_pipeline_run.  With d_qnodes and d_pair_nodes both null only the screen runs (the domain
 * sums and GMCR pairs need nothing else): d_valid is complete, d_counts holds only the zeros of
 * the pairs outside the GMCR.  No counterpart in the reference: it feeds the learned matching
 * order. */
int bms2_pipeline_trace(bms2_pipeline p, uint32_t max_graphs, uint64_t* d_counts,
                        uint64_t* d_valid, uint64_t* d_domsum, uint64_t* d_qpairs,
                        uint64_t* d_qnodes, uint32_t* d_pair_nodes, uint32_t* graphs_out);

/* ---------------------------------------------------------------- matching-order network
 * No counterpart in the reference (its join fixes identity order).  A multilayer perceptron
 * 32 -> 64 -> 64 -> 64 -> 10 with ReLU on the three hidden layers, one row per query node:
 * outputs [0] order score (higher = matched earlier), [1] predicted log2 of the DFS-node ratio
 * learned / identity order for the node's query, [2..9] logits of the useful refinement depth
 * 1..8.  Parameters are one f32 array: W1[32][64], b1[64], W2[64][64], b2[64], W3[64][64],
 * b3[64], W4[64][10], b4[10] (W[i][j] connects input i to output j).
 * Loss over a batch of rows r with targets t_r = {score, log2 ratio, depth index 0..7, weight w}:
 *   L = sum_r w_r [(y0 - t0)^2 + (y1 - t1)^2 + depth_weight * CE(softmax(y2..9), t2)] / sum_r w_r.
 * Training is f32 with Adam; inference is f32 or bf16 (weights and activations rounded to bf16
 * to nearest even, f32 accumulation). */
#define BMS2_MLP_IN 32
#define BMS2_MLP_HIDDEN 64
#define BMS2_MLP_OUT 10
#define BMS2_MLP_TARGETS 4

typedef struct bms2_adam {
  float lr, beta1, beta2, eps;
} bms2_adam;

/* Query-side input and output of the network, on the host (queries of at most 32 nodes).
 * bms2_query_static_features fills, per query node (rows in global node order, BMS2_MLP_IN floats
 * each), the feature columns that depend on the query's structure only -- 0-5 query size,
 * edges, cycle rank, largest degree, diameter, distinct labels; 8-10 degree, eccentricity,
 * closeness; 13-25 leaf neighbours, summed neighbour degrees, ring membership, wildcard and
 * element one-hot, position; 28 degree rank; 30 largest-degree flag -- and sets the others to 0.
 * bms2_score_orders writes order[32 g + d] for every query g: with use[g] != 0 the connected order
 * by descending scores[node] (the best node first, then always the best node adjacent to the
 * matched set, any remaining node when none is, ties to the lower id), identity otherwise.
 * No counterpart in the reference. */
int bms2_query_static_features(const bms2_host_csr* query, float* out);
int bms2_score_orders(const bms2_host_csr* query, const float* scores, const uint8_t* use,
                      uint8_t* order);

typedef struct bms2_mlp_s* bms2_mlp;
uint64_t bms2_mlp_num_params(void);
/* He-uniform weights from a splitmix64 stream of `seed`, zero biases, zero Adam state. */
int bms2_mlp_create(bms2_queue q, uint64_t seed, bms2_mlp* out);
int bms2_mlp_destroy(bms2_mlp m);
int bms2_mlp_get_params(bms2_mlp m, float* host, uint64_t n);
/* Replaces the parameters and resets the Adam state. */
int bms2_mlp_set_params(bms2_mlp m, const float* host, uint64_t n);
/* d_x: f32 [rows][32] (device); d_y: f32 [rows][10] (device). */
int bms2_mlp_forward(bms2_mlp m, const float* d_x, uint32_t rows, float* d_y, int bf16);
/* Loss and its gradient (grad_host: f32[num_params], nullable) at the current parameters;
 * d_t: f32 [rows][4] targets (device). */
int bms2_mlp_loss_grad(bms2_mlp m, const float* d_x, const float* d_t, uint32_t rows,
                       float depth_weight, float* loss, float* grad_host);
/* One Adam step on the batch; *loss (nullable) receives the loss before the step. */
int bms2_mlp_train_step(bms2_mlp m, const float* d_x, const float* d_t, uint32_t rows,
                        float depth_weight, const bms2_adam* opt, float* loss);

#ifdef __cplusplus
}
#endif
#endif /* BMS2_H */
