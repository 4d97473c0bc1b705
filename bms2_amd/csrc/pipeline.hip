// Fused batched pipeline: signatures -> filter -> K refinement steps -> GMCR -> join in one
// launch per resident data batch (the stage sequence of BMS2.cpp:131-234).
//
// One wave64 owns one pack of whole data graphs (<= 64 nodes, <= 8 graphs) for its lifetime:
//   1. node signature tuples (label, degree, path Bloom, ring counts, shell histograms) are
//      computed on chip, one node per lane, and kept in VGPRs;
//   2. every query is screened against the pack: its nodes' signature records come in through
//      scalar loads, each node's domain is one u64 ballot over the lanes, and a graph of the pack
//      drops out as soon as one query node has no candidate in it (GMCR validity, gmcr.hpp:54-88);
//   3. surviving (query, graph) pairs go to an LDS queue; the wave counts their embeddings with one
//      DFS per lane (isomorphism.hpp:291-472 semantics, standard backtracking), lanes refilling from
//      the queue as their pair finishes, so divergent DFS sizes do not idle the wave.
// Results: u64 count per (data graph, query), data-graph major, and GMCR membership bits.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <type_traits>
#include <vector>

#include "runtime.hpp"
#include "tuples.hpp"

namespace bms2 {

constexpr int kPackGraphs = 8;                 // graphs per pack
constexpr int kQueueSlots = kWave + kPackGraphs;  // LDS pair records

// Query node signature record, read with scalar loads (48 dwords).
struct alignas(16) QNode {
  uint32_t label;
  uint32_t deg;
  uint32_t level_mask;  // bit k: shell k+1 histogram non-zero
  uint32_t pad0;
  uint32_t path[8];
  uint32_t cyc[2];
  uint32_t pad1[2];
  // Per level: even / odd nibbles of the low and high histogram words, each nibble in its own
  // byte, plus 15 per byte: byte = q + 15, so (byte - d) has bit 4 set exactly when q > d.
  uint32_t hist[kMaxLevels][4];
};
static_assert(sizeof(QNode) == 192, "QNode layout");

// Per-query metadata (32 dwords).
struct alignas(16) QMeta {
  uint32_t node_begin;
  uint32_t nq;
  uint32_t valid;       // nq > 1 (gmcr.hpp:64)
  uint32_t pad;
  uint8_t forder[32];   // screening order of local query nodes (most selective first)
  uint8_t jorder[32];   // matching order: query node matched at each DFS depth
  uint32_t pad2[4];
};
static_assert(sizeof(QMeta) == 96, "QMeta layout");

struct FusedParams {
  const uint32_t* graph_offsets;
  const uint32_t* row_offsets;
  const uint32_t* column_indices;
  const uint8_t* labels;
  const uint32_t* pack_gbegin;  // [num_packs + 1]
  uint32_t num_packs;
  uint32_t Q;
  const QMeta* qmeta;
  const QNode* qnodes;
  const uint32_t* qadj;  // [Q][32] in matching order: bit i of word d = depth i adjacent to d
  uint64_t* counts;      // [D][Q]
  uint64_t* valid;       // [D][VW]
  uint64_t* stats;       // [4] or null
  uint32_t VW;
  uint32_t nq_cap;       // max query nodes, sizes the LDS records and stacks
  uint32_t find_first;
};

struct QueryPrepParams {
  const uint32_t* graph_offsets;
  const uint32_t* row_offsets;
  const uint32_t* column_indices;
  const uint8_t* labels;
  const uint32_t* pack_gbegin;
  uint32_t num_packs;
  QNode* qnodes;
};

// LDS layout of the fused kernel (per wave; one wave per workgroup).
struct FusedLds {
  uint64_t* adj;    // [64]
  uint64_t* m;      // [32] domain word per local query node of the current query
  uint8_t* lbl;     // [64]
  uint32_t* rec;    // [kQueueSlots][rec_words]  (aliases the Bloom scratch before the query loop)
  uint64_t* rstk;   // [nq_cap][64] remaining candidates per depth
  uint8_t* fstk;    // [nq_cap][64] matched data node per depth
  uint32_t rec_words;
};

__host__ __device__ inline uint32_t fused_rec_words(uint32_t nq_cap) { return 2 + 3 * nq_cap; }

__host__ __device__ inline size_t fused_lds_bytes(uint32_t nq_cap) {
  size_t b = 64 * 8 + 32 * 8 + 64;
  size_t rec = size_t(kQueueSlots) * fused_rec_words(nq_cap) * 4;
  b += std::max<size_t>(rec, 64 * 8 * 4);
  b += size_t(nq_cap) * 64 * 8 + size_t(nq_cap) * 64;
  return (b + 15) & ~size_t(15);
}

__device__ inline FusedLds carve(uint8_t* base, uint32_t nq_cap) {
  FusedLds l;
  l.adj = reinterpret_cast<uint64_t*>(base);
  base += 64 * 8;
  l.m = reinterpret_cast<uint64_t*>(base);
  base += 32 * 8;
  l.lbl = base;
  base += 64;
  l.rec = reinterpret_cast<uint32_t*>(base);
  l.rec_words = fused_rec_words(nq_cap);
  base += std::max<size_t>(size_t(kQueueSlots) * l.rec_words * 4, 64 * 8 * 4);
  l.rstk = reinterpret_cast<uint64_t*>(base);
  base += size_t(nq_cap) * 64 * 8;
  l.fstk = base;
  return l;
}

// ------------------------------------------------------------------ query signature records

template <int K>
__global__ __launch_bounds__(64) void query_prep_kernel(QueryPrepParams p) {
  __shared__ uint64_t adj[64];
  __shared__ uint8_t lbl[64];
  __shared__ uint32_t bloom[64 * 8];
  const uint32_t pack = blockIdx.x;
  const uint32_t gb = p.pack_gbegin[pack], ge = p.pack_gbegin[pack + 1];
  const uint32_t n0 = p.graph_offsets[gb], nn = p.graph_offsets[ge] - n0;
  Tuple t;
  compute_tuples<K>(n0, nn, p.row_offsets, p.column_indices, p.labels, PackScratch{adj, lbl, bloom},
                    t);
  const uint32_t lane = lane_id();
  if (lane >= nn) return;
  QNode r;
  r.label = t.label;
  r.deg = t.deg;
  r.level_mask = 0;
  r.pad0 = 0;
#pragma unroll
  for (int w = 0; w < 8; ++w) r.path[w] = t.path[w];
  r.cyc[0] = uint32_t(t.cycle);
  r.cyc[1] = uint32_t(t.cycle >> 32);
  r.pad1[0] = r.pad1[1] = 0;
#pragma unroll
  for (int k = 0; k < kMaxLevels; ++k) {
    const uint64_t h = t.hist[k];
    if (h) r.level_mask |= 1u << k;
    const uint32_t lo = uint32_t(h), hi = uint32_t(h >> 32);
    r.hist[k][0] = (lo & 0x0f0f0f0fu) + 0x0f0f0f0fu;
    r.hist[k][1] = ((lo >> 4) & 0x0f0f0f0fu) + 0x0f0f0f0fu;
    r.hist[k][2] = (hi & 0x0f0f0f0fu) + 0x0f0f0f0fu;
    r.hist[k][3] = ((hi >> 4) & 0x0f0f0f0fu) + 0x0f0f0f0fu;
  }
  p.qnodes[n0 + lane] = r;
}

// ------------------------------------------------------------------ fused screen + join kernel

template <int K>
__global__ __launch_bounds__(64) void fused_kernel(FusedParams p) {
  extern __shared__ __align__(16) uint8_t smem[];
  const FusedLds L = carve(smem, p.nq_cap);
  const uint32_t lane = lane_id();
  const uint32_t pack = blockIdx.x;
  const uint32_t gb = p.pack_gbegin[pack], ge = p.pack_gbegin[pack + 1];
  const uint32_t ng = ge - gb;
  const uint32_t n0 = p.graph_offsets[gb], nn = p.graph_offsets[ge] - n0;
  const uint64_t node_mask = low_mask(nn);

  // ---- 1. data node tuples (VGPR resident)
  Tuple t;
  compute_tuples<K>(n0, nn, p.row_offsets, p.column_indices, p.labels,
                    PackScratch{L.adj, L.lbl, L.rec}, t);
  const uint32_t d_label = lane < nn ? t.label : 0x100u;  // never equal to a u8 query label
  const uint32_t d_deg = t.deg;
  uint32_t d_pnot[8];
#pragma unroll
  for (int w = 0; w < 8; ++w) d_pnot[w] = ~t.path[w];
  // Ring counts with the top bit of each u16 set: (d | 0x8000) - q has bit 15 set iff d >= q.
  const uint32_t d_cyc0 = uint32_t(t.cycle) | 0x80008000u;
  const uint32_t d_cyc1 = uint32_t(t.cycle >> 32) | 0x80008000u;
  uint32_t d_hist[K > 0 ? K : 1][4];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint32_t lo = uint32_t(t.hist[k]), hi = uint32_t(t.hist[k] >> 32);
    d_hist[k][0] = lo & 0x0f0f0f0fu;
    d_hist[k][1] = (lo >> 4) & 0x0f0f0f0fu;
    d_hist[k][2] = hi & 0x0f0f0f0fu;
    d_hist[k][3] = (hi >> 4) & 0x0f0f0f0fu;
  }

  // graph lane masks of the pack (uniform)
  uint64_t gm[kPackGraphs];
#pragma unroll
  for (int j = 0; j < kPackGraphs; ++j) {
    gm[j] = 0;
    if (uint32_t(j) < ng) {
      const uint32_t s = p.graph_offsets[gb + j] - n0, e = p.graph_offsets[gb + j + 1] - n0;
      gm[j] = low_mask(e) & ~low_mask(s);
    }
  }

  // ---- join state (per lane) and queue state (uniform)
  int slot = -1;        // LDS record processed by this lane, -1 when idle
  int depth = 0;
  uint32_t jq = 0, jg = 0, jnq = 0;
  uint64_t used = 0, cnt = 0, nodes = 0;
  uint64_t st_matched = 0, st_emb = 0, st_nodes = 0;
  uint64_t free_lo = ~uint64_t(0), free_hi = low_mask(kQueueSlots - 64);
  uint64_t pend_lo = 0, pend_hi = 0;
  uint64_t st_pairs = 0;
  uint64_t vbits = 0;  // lane j < ng: GMCR bits of graph j for the current block of 64 queries
  const uint32_t RW = L.rec_words;
  const uint32_t ncap = p.nq_cap;

  auto rec_ptr = [&](int s) { return L.rec + size_t(s) * RW; };

  // One DFS step for every busy lane; returns nothing, updates lane state.
  auto join_phase = [&](bool drain) {
    while (true) {
      // hand pending records to idle lanes
      uint64_t idle = ballot(slot < 0);
      bool fresh = false;
      while (idle && (pend_lo | pend_hi)) {
        const int l = ctz64(idle);
        idle &= idle - 1;
        int s;
        if (pend_lo) {
          s = ctz64(pend_lo);
          pend_lo &= pend_lo - 1;
        } else {
          s = 64 + ctz64(pend_hi);
          pend_hi &= pend_hi - 1;
        }
        if (int(lane) == l) {
          slot = s;
          fresh = true;
        }
      }
      if (fresh) {
        const uint32_t* r = rec_ptr(slot);
        jq = r[0];
        jg = r[1] & 0xffu;
        jnq = r[1] >> 8;
        depth = 0;
        used = 0;
        cnt = 0;
        nodes = 0;
        L.rstk[lane] = *reinterpret_cast<const uint64_t*>(r + 2);  // domain of depth 0
      }
      const uint64_t busy = ballot(slot >= 0);
      if (!busy) break;
      if (!drain && !(pend_lo | pend_hi) && busy != ~uint64_t(0)) break;
      bool finished = false;
      if (slot >= 0) {
        const uint32_t* r = rec_ptr(slot);
        const uint64_t* dom = reinterpret_cast<const uint64_t*>(r + 2);
        const uint32_t* qa = r + 2 + 2 * ncap;
        uint64_t R = L.rstk[depth * 64 + lane];
        if (R == 0) {
          if (depth == 0) {
            finished = true;
          } else {
            --depth;
            used &= ~(uint64_t(1) << L.fstk[depth * 64 + lane]);
          }
        } else {
          const int v = ctz64(R);
          L.rstk[depth * 64 + lane] = R & (R - 1);
          L.fstk[depth * 64 + lane] = uint8_t(v);
          ++nodes;
          const int nd = depth + 1;
          // candidates of depth nd: domain, unused, and query adjacency == data adjacency to
          // every matched node (induced, isomorphism.hpp:268-288)
          const uint64_t vbit = uint64_t(1) << v;
          uint64_t Rn = dom[nd] & ~(used | vbit);
          const uint32_t am = qa[nd];
          for (int i = 0; i < nd && Rn; ++i) {
            const uint64_t a = L.adj[L.fstk[i * 64 + lane]];
            Rn &= ((am >> i) & 1) ? a : ~a;
          }
          if (uint32_t(nd) == jnq - 1) {
            cnt += uint64_t(popc64(Rn));
            if (p.find_first && cnt) finished = true;
          } else {
            used |= vbit;
            depth = nd;
            L.rstk[nd * 64 + lane] = Rn;
          }
        }
      }
      if (finished) {
        const uint64_t c = p.find_first ? (cnt ? 1 : 0) : cnt;
        p.counts[uint64_t(gb + jg) * p.Q + jq] = c;
        st_matched += c ? 1 : 0;
        st_emb += c;
        st_nodes += nodes;
      }
      uint64_t fin = ballot(finished);
      while (fin) {
        const int l = ctz64(fin);
        fin &= fin - 1;
        const int s = __builtin_amdgcn_readlane(slot, l);
        if (s < 64)
          free_lo |= uint64_t(1) << s;
        else
          free_hi |= uint64_t(1) << (s - 64);
      }
      if (finished) slot = -1;
    }
  };

  // ---- 2. screen every query against the pack
  for (uint32_t q = 0; q < p.Q; ++q) {
    const QMeta* meta = p.qmeta + q;
    const uint32_t nq = meta->nq;
    uint32_t alive = meta->valid ? (ng >= 32 ? ~0u : ((1u << ng) - 1)) : 0u;
    uint64_t lanes_alive = node_mask;
    if (alive) {
      const QNode* qbase = p.qnodes + meta->node_begin;
      for (uint32_t ti = 0; ti < nq; ++ti) {
        const uint32_t i = meta->forder[ti];
        const QNode* qn = qbase + i;
        const uint32_t ql = qn->label;
        uint64_t m = ballot((d_label == ql || ql == kWildcard) && d_deg >= qn->deg) & lanes_alive;
        if (m) {
          uint32_t acc = 0;
#pragma unroll
          for (int w = 0; w < 8; ++w) acc |= qn->path[w] & d_pnot[w];
          m &= ballot(acc == 0);
        }
        if (m) {
          const uint32_t c0 = d_cyc0 - qn->cyc[0], c1 = d_cyc1 - qn->cyc[1];
          m &= ballot(((c0 & c1) & 0x80008000u) == 0x80008000u);
        }
        if (K > 0 && m) {
          const uint32_t lm = qn->level_mask;
          uint32_t fail = 0;
#pragma unroll
          for (int k = 0; k < K; ++k) {
            if ((lm >> k) & 1) {
              fail |= (qn->hist[k][0] - d_hist[k][0]) | (qn->hist[k][1] - d_hist[k][1]) |
                      (qn->hist[k][2] - d_hist[k][2]) | (qn->hist[k][3] - d_hist[k][3]);
            }
          }
          m &= ballot((fail & 0x10101010u) == 0);
        }
        if (lane == 0) L.m[i] = m;
#pragma unroll
        for (int j = 0; j < kPackGraphs; ++j) {
          if (((alive >> j) & 1) && !(m & gm[j])) {
            alive &= ~(1u << j);
            lanes_alive &= ~gm[j];
          }
        }
        if (!alive) break;
      }
    }
    __syncthreads();  // L.m visible to every lane
    // pairs outside the GMCR: count 0
    if (lane < ng && !((alive >> lane) & 1)) p.counts[uint64_t(gb + lane) * p.Q + q] = 0;
    if (lane < ng) vbits |= uint64_t((alive >> lane) & 1) << (q & 63);
    if ((q & 63) == 63 || q + 1 == p.Q) {
      if (lane < ng) p.valid[uint64_t(gb + lane) * p.VW + (q >> 6)] = vbits;
      vbits = 0;
    }
    // GMCR pairs -> LDS records (domains and adjacency in matching order)
    while (alive) {
      const int j = __builtin_ctz(alive);
      alive &= alive - 1;
      int s;
      if (free_lo) {
        s = ctz64(free_lo);
        free_lo &= free_lo - 1;
        pend_lo |= uint64_t(1) << s;
      } else {
        s = 64 + ctz64(free_hi);
        free_hi &= free_hi - 1;
        pend_hi |= uint64_t(1) << (s - 64);
      }
      ++st_pairs;
      uint32_t* r = rec_ptr(s);
      uint64_t gmj = 0;
#pragma unroll
      for (int jj = 0; jj < kPackGraphs; ++jj)
        if (jj == j) gmj = gm[jj];
      if (lane == 0) {
        r[0] = q;
        r[1] = uint32_t(j) | (nq << 8);
      }
      if (lane < nq) {
        const uint32_t node = meta->jorder[lane];
        reinterpret_cast<uint64_t*>(r + 2)[lane] = L.m[node] & gmj;
        r[2 + 2 * ncap + lane] = p.qadj[size_t(q) * 32 + lane];
      }
    }
    __syncthreads();
    const uint64_t idle = ballot(slot < 0);
    const uint32_t npend = uint32_t(popc64(pend_lo) + popc64(pend_hi));
    if (npend && npend >= uint32_t(popc64(idle))) join_phase(false);
  }
  join_phase(true);

  // ---- stats
  if (p.stats) {
    for (int off = 32; off > 0; off >>= 1) {
      st_matched += __shfl_xor(st_matched, off);
      st_emb += __shfl_xor(st_emb, off);
      st_nodes += __shfl_xor(st_nodes, off);
    }
    if (lane == 0) {
      atomicAdd(reinterpret_cast<unsigned long long*>(p.stats + 0), (unsigned long long)st_pairs);
      atomicAdd(reinterpret_cast<unsigned long long*>(p.stats + 1), (unsigned long long)st_matched);
      atomicAdd(reinterpret_cast<unsigned long long*>(p.stats + 2), (unsigned long long)st_emb);
      atomicAdd(reinterpret_cast<unsigned long long*>(p.stats + 3), (unsigned long long)st_nodes);
    }
  }
}

// ------------------------------------------------------------------ host side

// Greedy packing of consecutive graphs: at most 64 nodes and kPackGraphs graphs per pack.
std::vector<uint32_t> build_packs(const std::vector<uint32_t>& graph_offsets) {
  const uint32_t G = uint32_t(graph_offsets.size() - 1);
  std::vector<uint32_t> begins;
  uint32_t g = 0;
  while (g < G) {
    begins.push_back(g);
    uint32_t nodes = 0, cnt = 0;
    while (g < G && cnt < uint32_t(kPackGraphs)) {
      const uint32_t n = graph_offsets[g + 1] - graph_offsets[g];
      if (cnt > 0 && nodes + n > 64) break;
      nodes += n;
      ++cnt;
      ++g;
    }
  }
  begins.push_back(G);
  return begins;
}

template <class F>
int dispatch_levels(uint32_t K, F&& f) {
  switch (K) {
    case 0: return f(std::integral_constant<int, 0>{});
    case 1: return f(std::integral_constant<int, 1>{});
    case 2: return f(std::integral_constant<int, 2>{});
    case 3: return f(std::integral_constant<int, 3>{});
    case 4: return f(std::integral_constant<int, 4>{});
    case 5: return f(std::integral_constant<int, 5>{});
    case 6: return f(std::integral_constant<int, 6>{});
    case 7: return f(std::integral_constant<int, 7>{});
    case 8: return f(std::integral_constant<int, 8>{});
  }
  return set_error(BMS2_ERR_UNSUPPORTED, "refinement steps %u > %d", K, kMaxLevels);
}

}  // namespace bms2

using namespace bms2;

struct bms2_pipeline_s {
  bms2_queue q;
  bms2_device_csr data, query;
  bms2_pipeline_options opt;
  uint32_t num_packs = 0, num_qpacks = 0, nq_cap = 0;
  uint32_t* d_packs = nullptr;
  uint32_t* d_qpacks = nullptr;
  QMeta* d_qmeta = nullptr;
  QNode* d_qnodes = nullptr;
  uint32_t* d_qadj = nullptr;
};

namespace {

void pipeline_free(bms2_pipeline p) {
  if (!p) return;
  (void)hipFree(p->d_packs);
  (void)hipFree(p->d_qpacks);
  (void)hipFree(p->d_qmeta);
  (void)hipFree(p->d_qnodes);
  (void)hipFree(p->d_qadj);
  delete p;
}

}  // namespace

extern "C" int bms2_pipeline_create(bms2_queue q, const bms2_device_csr* query,
                                    const bms2_device_csr* data, const bms2_pipeline_options* opt,
                                    bms2_pipeline* out) {
  if (int rc = use_device(q)) return rc;
  if (!query || !data || !opt || !out) return set_error(BMS2_ERR_INVALID, "null argument");
  if (opt->refinement_steps > uint32_t(kMaxLevels))
    return set_error(BMS2_ERR_UNSUPPORTED, "refinement steps %u > %d", opt->refinement_steps,
                     kMaxLevels);
  if (data->max_graph_nodes > 64)
    return set_error(BMS2_ERR_UNSUPPORTED,
                     "fused pipeline handles data graphs of up to 64 nodes (largest: %u); use the "
                     "staged entry points",
                     data->max_graph_nodes);
  if (data->max_degree > 27 || query->max_degree > 27)
    return set_error(BMS2_ERR_UNSUPPORTED,
                     "node degree above 27: the reference's 128-entry ring-count stack could "
                     "overflow (signature.hpp:638,697), which the fused ring counts do not model");
  if (query->max_graph_nodes > BMS2_MAX_QUERY_NODES)
    return set_error(BMS2_ERR_UNSUPPORTED, "query with %u nodes (limit %u)",
                     query->max_graph_nodes, BMS2_MAX_QUERY_NODES);
  auto* p = new bms2_pipeline_s{};
  p->q = q;
  p->data = *data;
  p->query = *query;
  p->opt = *opt;
  p->nq_cap = query->max_graph_nodes <= 16 ? 16 : 32;
  hipStream_t s = q->stream;
  int rc = BMS2_OK;
  auto fail = [&](int code) {
    pipeline_free(p);
    return code;
  };
  // data packs (plan built once per resident batch)
  std::vector<uint32_t> go(data->num_graphs + 1);
  if (hipMemcpy(go.data(), data->graph_offsets, 4ull * go.size(), hipMemcpyDeviceToHost) !=
      hipSuccess)
    return fail(set_error(BMS2_ERR_HIP, "copy of data graph offsets failed"));
  std::vector<uint32_t> packs = build_packs(go);
  p->num_packs = uint32_t(packs.size() - 1);
  if ((rc = device_alloc((void**)&p->d_packs, 4ull * packs.size()))) return fail(rc);
  BMS2_HIP(hipMemcpyAsync(p->d_packs, packs.data(), 4ull * packs.size(), hipMemcpyHostToDevice, s));

  // queries: host copy of the small query batch for orders and adjacency
  const uint32_t Q = query->num_graphs;
  std::vector<uint32_t> qgo(Q + 1), qro(query->total_nodes + 1), qci(query->total_edges);
  std::vector<uint8_t> qlab(query->total_nodes);
  if (hipMemcpy(qgo.data(), query->graph_offsets, 4ull * qgo.size(), hipMemcpyDeviceToHost) ||
      hipMemcpy(qro.data(), query->row_offsets, 4ull * qro.size(), hipMemcpyDeviceToHost) ||
      (qci.size() && hipMemcpy(qci.data(), query->column_indices, 4ull * qci.size(),
                               hipMemcpyDeviceToHost)) ||
      (qlab.size() && hipMemcpy(qlab.data(), query->node_labels, qlab.size(), hipMemcpyDeviceToHost)))
    return fail(set_error(BMS2_ERR_HIP, "copy of the query batch failed"));
  std::vector<uint32_t> label_freq(256, 0);
  for (uint8_t l : qlab) label_freq[l]++;
  std::vector<QMeta> meta(Q);
  std::vector<uint32_t> qadj(size_t(Q) * 32, 0);
  for (uint32_t g = 0; g < Q; ++g) {
    QMeta& m = meta[g];
    std::memset(&m, 0, sizeof(m));
    const uint32_t b = qgo[g], n = qgo[g + 1] - qgo[g];
    m.node_begin = b;
    m.nq = n;
    m.valid = n > 1 ? 1 : 0;
    // screening order: rarest label first, then highest degree (only affects speed)
    std::vector<uint32_t> ord(n);
    std::iota(ord.begin(), ord.end(), 0u);
    std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t c) {
      const uint32_t fa = qlab[b + a] == 255 ? ~0u : label_freq[qlab[b + a]];
      const uint32_t fc = qlab[b + c] == 255 ? ~0u : label_freq[qlab[b + c]];
      if (fa != fc) return fa < fc;
      return qro[b + a + 1] - qro[b + a] > qro[b + c + 1] - qro[b + c];
    });
    for (uint32_t i = 0; i < n; ++i) {
      m.forder[i] = uint8_t(ord[i]);
      m.jorder[i] = uint8_t(i);  // identity matching order (isomorphism.hpp:371-447)
    }
    for (uint32_t d = 0; d < n; ++d) {
      const uint32_t u = b + m.jorder[d];
      uint32_t w = 0;
      for (uint32_t i = 0; i < d; ++i) {
        const uint32_t v = b + m.jorder[i];
        for (uint32_t k = qro[u]; k < qro[u + 1]; ++k)
          if (qci[k] == v) {
            w |= 1u << i;
            break;
          }
      }
      qadj[size_t(g) * 32 + d] = w;
    }
  }
  std::vector<uint32_t> qpacks = build_packs(qgo);
  p->num_qpacks = uint32_t(qpacks.size() - 1);
  if ((rc = device_alloc((void**)&p->d_qpacks, 4ull * qpacks.size())) ||
      (rc = device_alloc((void**)&p->d_qmeta, sizeof(QMeta) * std::max<size_t>(Q, 1))) ||
      (rc = device_alloc((void**)&p->d_qnodes,
                         sizeof(QNode) * std::max<uint64_t>(query->total_nodes, 1))) ||
      (rc = device_alloc((void**)&p->d_qadj, 4ull * std::max<size_t>(qadj.size(), 1))))
    return fail(rc);
  BMS2_HIP(hipMemcpyAsync(p->d_qpacks, qpacks.data(), 4ull * qpacks.size(), hipMemcpyHostToDevice, s));
  if (Q) {
    BMS2_HIP(hipMemcpyAsync(p->d_qmeta, meta.data(), sizeof(QMeta) * Q, hipMemcpyHostToDevice, s));
    BMS2_HIP(hipMemcpyAsync(p->d_qadj, qadj.data(), 4ull * qadj.size(), hipMemcpyHostToDevice, s));
  }
  BMS2_HIP(hipStreamSynchronize(s));
  *out = p;
  return BMS2_OK;
}

extern "C" int bms2_pipeline_destroy(bms2_pipeline p) {
  if (!p) return BMS2_OK;
  if (int rc = use_device(p->q)) return rc;
  (void)hipStreamSynchronize(p->q->stream);
  pipeline_free(p);
  return BMS2_OK;
}

extern "C" int bms2_pipeline_run(bms2_pipeline p, uint64_t* d_counts, uint64_t* d_valid,
                                 uint64_t* d_stats, bms2_event* ev, bms2_event* kernel_ev) {
  if (!p) return set_error(BMS2_ERR_INVALID, "pipeline is null");
  if (int rc = use_device(p->q)) return rc;
  if (!d_counts || !d_valid) return set_error(BMS2_ERR_INVALID, "null output buffer");
  hipStream_t s = p->q->stream;
  EventScope es{p->q, ev};
  if (int rc = es.begin()) return rc;
  const uint32_t Q = p->query.num_graphs;
  if (d_stats) BMS2_HIP(hipMemsetAsync(d_stats, 0, 8 * 8, s));
  if (p->num_packs == 0 || Q == 0) {
    BMS2_HIP(hipMemsetAsync(d_valid, 0, 8ull * p->data.num_graphs * ((Q + 63) / 64), s));
    return es.end();
  }
  QueryPrepParams qp{p->query.graph_offsets, p->query.row_offsets, p->query.column_indices,
                     p->query.node_labels, p->d_qpacks, p->num_qpacks, p->d_qnodes};
  FusedParams fp{};
  fp.graph_offsets = p->data.graph_offsets;
  fp.row_offsets = p->data.row_offsets;
  fp.column_indices = p->data.column_indices;
  fp.labels = p->data.node_labels;
  fp.pack_gbegin = p->d_packs;
  fp.num_packs = p->num_packs;
  fp.Q = Q;
  fp.qmeta = p->d_qmeta;
  fp.qnodes = p->d_qnodes;
  fp.qadj = p->d_qadj;
  fp.counts = d_counts;
  fp.valid = d_valid;
  fp.stats = d_stats;
  fp.VW = (Q + 63) / 64;
  fp.nq_cap = p->nq_cap;
  fp.find_first = p->opt.find_first ? 1 : 0;
  const size_t lds = fused_lds_bytes(p->nq_cap);
  int rc = dispatch_levels(p->opt.refinement_steps, [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    if (p->num_qpacks)
      hipLaunchKernelGGL(query_prep_kernel<K>, dim3(p->num_qpacks), dim3(64), 0, s, qp);
    BMS2_CHECK_LAUNCH("query_prep_kernel");
    EventScope ks{p->q, kernel_ev};
    if (int krc = ks.begin()) return krc;
    hipLaunchKernelGGL(fused_kernel<K>, dim3(p->num_packs), dim3(64), lds, s, fp);
    BMS2_CHECK_LAUNCH("fused_kernel");
    return ks.end();
  });
  if (rc) return rc;
  return es.end();
}
