/*
 * oracle.h -- CPU restatement of the reference's batched subgraph-isomorphism path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load this library, and only as the checker (or the timed CPU baseline); the product
 * (bms2_amd / libbms2.so) never links or calls it.
 *
 * Each function follows the reference file:line it names (paths relative to AnySparse/BMS2).
 * Where the reference is undefined or buggy the oracle takes the evident intent:
 *   - DFS backtracking restores exactly what it set (isomorphism.hpp:418-447 leaks visited bits);
 *   - BFS / visited sets are exact for any graph size (utils.hpp:95-97 truncates ids to 8 bits
 *     and shifts past 64);
 *   - reachability state starts at zero (signature.hpp:162-166 leaves it uninitialised);
 *   - range masks are exact (candidates.hpp:94,140 shift by 32 when end % 32 == 0);
 *   - queries of up to 32 nodes (isomorphism.hpp:248 sizes arrays for 30);
 *   - GMCR segments list query ids in ascending order (gmcr.hpp:120-126 uses atomics).
 * Everything else -- 4-bit saturating histograms of labels < 16, the Bloom hash chain, the
 * literal cycle DFS with its 128-entry stack and 64-node cutoff, induced matching, single-node
 * queries never valid -- is reproduced as written.
 */
#ifndef BMS2_ORACLE_H
#define BMS2_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Packed batched CSR (layout of DeviceBatchedCSRGraph, graph.hpp:58-66). */
typedef struct orc_csr {
  uint32_t num_graphs;
  uint32_t reserved;
  uint64_t total_nodes;
  uint64_t total_edges;
  const uint32_t* graph_offsets;
  const uint32_t* row_offsets;
  const uint32_t* column_indices;
  const uint8_t* node_labels;
  const uint8_t* edge_labels;
} orc_csr;

/* MurmurHash3 fmix32 (signature.hpp:393-400) and boost hash_combine (:402-404). */
uint32_t orc_hash32(uint32_t h);
uint32_t orc_combine_hash(uint32_t h1, uint32_t h2);

/* generateCSRSignatures (signature.hpp:124-151): 4-bit saturating counts of neighbour labels. */
void orc_label_signatures(const orc_csr* g, uint64_t* sig);
/* generateCSRPathSignatures (signature.hpp:472-519); out is [node][4]. */
void orc_path_signatures(const orc_csr* g, uint64_t* out);
/* generateCSRCycleSignatures (signature.hpp:591-719); out is one u64 per node holding the u16
 * fields {5 same, 5 diff, 6 same, 6 diff} from the low end. */
void orc_cycle_signatures(const orc_csr* g, uint64_t* out);
/* refineCSRSignatures<PowerGraph> (signature.hpp:336-389); reach is [node][4] (256 bits). */
void orc_refine_signatures(const orc_csr* g, uint32_t view_size, uint64_t* sig, uint64_t* reach);

/* filterCandidates (isomorphism.hpp:30-104): OR-s bits into cand ([Qn][W] u64 words). */
void orc_filter(const orc_csr* q, const orc_csr* d, const uint64_t* qpath, const uint64_t* dpath,
                const uint64_t* qcycle, const uint64_t* dcycle, uint64_t* cand, uint64_t W);
/* refineCandidates (isomorphism.hpp:106-203). */
void orc_refine_candidates(const orc_csr* q, const orc_csr* d, const uint64_t* qsig,
                           const uint64_t* dsig, uint64_t* cand, uint64_t W);
/* generateGMCR (gmcr.hpp:35-142).  Returns the number of pairs; fills indices up to cap. */
uint64_t orc_gmcr(const orc_csr* q, const orc_csr* d, const uint64_t* cand, uint64_t W,
                  uint32_t* offsets, uint32_t* indices, uint64_t cap);
/* joinCandidates (isomorphism.hpp:291-472), standard backtracking.  `order` (nullable) holds,
 * per global query node slot, the local query node matched at that depth (a permutation per
 * query); null means identity order.  counts[t] per GMCR entry; dfs_nodes[t] (nullable) counts
 * partial embeddings of every length below n_q.  Returns the total. */
uint64_t orc_join(const orc_csr* q, const orc_csr* d, const uint64_t* cand, uint64_t W,
                  const uint32_t* offsets, const uint32_t* indices, int find_first,
                  const uint32_t* order, uint64_t* counts, uint64_t* dfs_nodes);
/* orc_join with a symmetry-breaking constraint per query (not a reference behaviour: the checker of
 * the fused pipeline's large-query join): at a depth in sym_vmask[qi] a candidate must exceed the
 * data node matched at depth sym_du[qi] (0xff: none).  counts are the constrained counts. */
uint64_t orc_join_sym(const orc_csr* q, const orc_csr* d, const uint64_t* cand, uint64_t W,
                      const uint32_t* offsets, const uint32_t* indices, int find_first,
                      const uint32_t* order, const uint8_t* sym_du, const uint32_t* sym_vmask,
                      uint64_t* counts, uint64_t* dfs_nodes);
/* joinPartialCandidates2 (isomorphism.hpp:1021-1178), standard backtracking: every embedding of
 * query 0's first `end` nodes in every data graph, as records of end + 1 ints {mapped global data
 * node ids, data graph id} in data graph order then DFS order; writes at most cap records (out
 * nullable) and returns how many there are. */
uint64_t orc_join_partial(const orc_csr* q, const orc_csr* d, const uint64_t* cand, uint64_t W,
                          uint32_t end, int32_t* out, uint64_t cap);
/* joinWithPartialMatches (isomorphism.hpp:756-977), standard backtracking: counts[t] (per GMCR
 * entry, overwritten) = over the partial matches of entry t's data graph, 1 each when n_q == end,
 * their extensions to all of q when n_q > end; find_first keeps [count > 0].  Returns the total. */
uint64_t orc_join_with_partial(const orc_csr* q, const orc_csr* d, const uint64_t* cand,
                               uint64_t W, const uint32_t* offsets, const uint32_t* indices,
                               uint32_t end, const int32_t* partial, uint64_t num_partial,
                               int find_first, uint64_t* counts);

/* Fused per-data-graph restatement of the stage sequence BMS2.cpp:131-234 with K refinement
 * steps, over data graphs [g_begin, g_end), parallel over data graphs (OpenMP, `threads`, 0 =
 * default).  counts: [(g_end-g_begin)][Q] u64; valid: [(g_end-g_begin)][ceil(Q/64)] bit rows;
 * stats (nullable) u64[4] {GMCR pairs, matched pairs, embeddings, DFS nodes}.  This is the timed
 * CPU baseline; it must equal the staged functions above. */
int orc_pipeline(const orc_csr* q, const orc_csr* d, uint32_t K, int find_first,
                 uint32_t g_begin, uint32_t g_end, uint64_t* counts, uint64_t* valid,
                 uint64_t* stats, int threads);

#ifdef __cplusplus
}
#endif
#endif
